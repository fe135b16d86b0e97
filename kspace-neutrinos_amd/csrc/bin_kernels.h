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
inline long long build_thresholds(int dims, int nrbins, std::vector<int> &thr)
{
    const double binsperunit = (nrbins - 1) / log(sqrt(3) * dims / 2.0);
    const long long kmax2 = 3LL * (dims / 2) * (dims / 2);
    thr.assign(nrbins + 1, 0x7fffffff);
    thr[0] = 0;
    int prev = 0;
    for (long long k2 = 1; k2 <= kmax2; k2++) {
        const int b = (int)floor(binsperunit * log(sqrt((double)k2)));
        if (b < prev || b >= nrbins) return k2;
        while (prev < b) thr[++prev] = (int)k2;
    }
    return 0;
}

// Per-axis inverse CiC window by storage index (powerspectrum.c:8-12).
inline void build_invwindow(int dims, std::vector<double> &iw)
{
    iw.resize(dims);
    for (int n = 0; n < dims; n++) {
        const int kx = kval(n, dims);
        iw[n] = kx ? M_PI * kx / (dims * sin(M_PI * kx / (double)dims)) : 1.0;
    }
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

template <int WAVES>
__host__ __device__ constexpr size_t smem_bytes(int dims, int nrbins)
{
    return sizeof(double) * ((size_t)WAVES * nrbins + dims) + sizeof(int) * (nrbins + 1);
}

template <int WAVES>
__device__ __forceinline__ void load_tables(const BinArgs &a, double *&hist, double *&iw, int *&thr)
{
    extern __shared__ double smem[];
    hist = smem;
    iw = hist + (size_t)WAVES * a.nrbins;
    thr = (int *)(iw + a.dims);
    for (int t = threadIdx.x; t < WAVES * a.nrbins; t += 64 * WAVES) hist[t] = 0.0;
    for (int t = threadIdx.x; t < a.dims; t += 64 * WAVES) iw[t] = a.iw[t];
    for (int t = threadIdx.x; t <= a.nrbins; t += 64 * WAVES) thr[t] = a.thr[t];
    __syncthreads();
}

template <int WAVES>
__device__ __forceinline__ void store_partial(const BinArgs &a, const double *hist)
{
    __syncthreads();
    for (int b = threadIdx.x; b < a.nrbins; b += 64 * WAVES) {
        double s = hist[b];
        for (int w = 1; w < WAVES; w++) s += hist[(size_t)w * a.nrbins + b];
        a.partial[(size_t)blockIdx.x * a.nrbins + b] = s;
    }
}

// Each lane streams RUN consecutive modes of its strip (stride RUN*16 B across
// lanes, so a wave's RUN loads cover one contiguous 1 KB*RUN span); bins are
// non-decreasing along a row, so equal-bin runs are summed in registers and
// flushed to the wave's LDS histogram once per run.
template <int WAVES, int RUN>
__global__ __launch_bounds__(64 * WAVES) void k_bin_power_runs(BinArgs a)
{
    double *hist, *iw;
    int *thr;
    load_tables<WAVES>(a, hist, iw, thr);
    constexpr int B = 64 * WAVES;
    const int rowlen = a.dims / 2 + 1, half = a.dims / 2;
    const long long r0 = a.rows * blockIdx.x / gridDim.x, r1 = a.rows * (blockIdx.x + 1) / gridDim.x;
    const long long m0 = r0 * rowlen, m1 = r1 * rowlen;
    double *my = hist + (size_t)(threadIdx.x >> 6) * a.nrbins;
    Pos p = pos_of((long long)threadIdx.x * RUN, r0, rowlen, a.dims);
    for (long long m = m0 + (long long)threadIdx.x * RUN; m < m1; m += (long long)B * RUN) {
        double2 f[RUN];
#pragma unroll
        for (int r = 0; r < RUN; r++) f[r] = (m + r < m1) ? a.F[m + r] : make_double2(0.0, 0.0);
        Pos q = p;
        int cur = -1;
        double acc = 0.0;
#pragma unroll
        for (int r = 0; r < RUN; r++) {
            if (m + r >= m1) break;
            if (r) advance(q, 1, rowlen, a.dims);
            const int kx = kval(a.startslab + q.i, a.dims), ky = kval(q.j, a.dims);
            const int k2 = kx * kx + ky * ky + q.z * q.z;
            if (k2 == 0) continue;
            const int b = find_bin(k2, a.half_bpu, thr, a.nrbins);
            const double v = mode_weight(f[r], __dmul_rn(iw[a.startslab + q.i], iw[q.j]), iw[q.z], q.z, half);
            if (b != cur) {
                if (cur >= 0) atomicAdd(&my[cur], acc);
                cur = b;
                acc = v;
            } else {
                acc += v;
            }
        }
        if (cur >= 0) atomicAdd(&my[cur], acc);
        advance(p, B * RUN, rowlen, a.dims);
    }
    store_partial<WAVES>(a, hist);
}

// Coalesced variant: lane l of the block takes modes l, l+B, l+2B, ... with
// UNROLL loads in flight.
template <int WAVES, int UNROLL>
__global__ __launch_bounds__(64 * WAVES) void k_bin_power_coalesced(BinArgs a)
{
    double *hist, *iw;
    int *thr;
    load_tables<WAVES>(a, hist, iw, thr);
    constexpr int B = 64 * WAVES;
    const int rowlen = a.dims / 2 + 1, half = a.dims / 2;
    const long long r0 = a.rows * blockIdx.x / gridDim.x, r1 = a.rows * (blockIdx.x + 1) / gridDim.x;
    const long long m1 = r1 * rowlen;
    double *my = hist + (size_t)(threadIdx.x >> 6) * a.nrbins;
    Pos p = pos_of(threadIdx.x, r0, rowlen, a.dims);
    for (long long m = r0 * rowlen + threadIdx.x; m < m1; m += (long long)B * UNROLL) {
        double2 f[UNROLL];
        Pos q[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            q[u] = p;
            const long long mu = m + (long long)B * u;
            f[u] = mu < m1 ? a.F[mu] : make_double2(0.0, 0.0);
            advance(p, B, rowlen, a.dims);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            if (m + (long long)B * u >= m1) break;
            const int kx = kval(a.startslab + q[u].i, a.dims), ky = kval(q[u].j, a.dims), kz = q[u].z;
            const int k2 = kx * kx + ky * ky + kz * kz;
            if (k2 == 0) continue;
            const int b = find_bin(k2, a.half_bpu, thr, a.nrbins);
            atomicAdd(&my[b], mode_weight(f[u], __dmul_rn(iw[a.startslab + q[u].i], iw[q[u].j]), iw[kz], kz, half));
        }
    }
    store_partial<WAVES>(a, hist);
}

// Software-pipelined runs: each lane bins RUN consecutive modes of the
// flattened grid while its next RUN modes are in flight.  Along a row |k|^2
// grows with z, so after the first mode of a run the bin is found
// incrementally: the next threshold is kept in a register and at most one
// bin step is taken before falling back to the full search (near the origin a
// unit step in z crosses many log bins).  Equal-bin runs are summed in
// registers and flushed to the wave's LDS histogram.
template <int WAVES, int RUN>
__global__ __launch_bounds__(64 * WAVES) void k_bin_power_runs_pf(BinArgs a)
{
    double *hist, *iw;
    int *thr;
    load_tables<WAVES>(a, hist, iw, thr);
    constexpr int B = 64 * WAVES;
    const int rowlen = a.dims / 2 + 1, half = a.dims / 2;
    const long long r0 = a.rows * blockIdx.x / gridDim.x, r1 = a.rows * (blockIdx.x + 1) / gridDim.x;
    const long long m0 = r0 * rowlen, m1 = r1 * rowlen;
    const long long S = (long long)B * RUN;
    double *my = hist + (size_t)(threadIdx.x >> 6) * a.nrbins;
    Pos p = pos_of((long long)threadIdx.x * RUN, r0, rowlen, a.dims);
    long long m = m0 + (long long)threadIdx.x * RUN;
    double2 cur[RUN];
#pragma unroll
    for (int r = 0; r < RUN; r++) cur[r] = (m + r < m1) ? a.F[m + r] : make_double2(0.0, 0.0);
    for (; m < m1; m += S) {
        double2 nxt[RUN];
#pragma unroll
        for (int r = 0; r < RUN; r++) nxt[r] = (m + S + r < m1) ? a.F[m + S + r] : make_double2(0.0, 0.0);
        Pos q = p;
        int curb = -1, b = -1, bnext = 0;
        double acc = 0.0;
#pragma unroll
        for (int r = 0; r < RUN; r++) {
            if (m + r >= m1) break;
            if (r) advance(q, 1, rowlen, a.dims);
            const int kx = kval(a.startslab + q.i, a.dims), ky = kval(q.j, a.dims);
            const int k2 = kx * kx + ky * ky + q.z * q.z;
            if (k2 == 0) continue;
            if (b < 0 || q.z == 0) {  // run start or a new row: full search
                b = find_bin(k2, a.half_bpu, thr, a.nrbins);
                bnext = thr[b + 1];
            } else if (k2 >= bnext) {
                const int t2 = thr[b + 2];  // here b + 1 <= nrbins - 1, so b + 2 <= nrbins
                if (k2 < t2) {
                    b++;
                    bnext = t2;
                } else {
                    b = find_bin(k2, a.half_bpu, thr, a.nrbins);
                    bnext = thr[b + 1];
                }
            }
            const double v = mode_weight(cur[r], __dmul_rn(iw[a.startslab + q.i], iw[q.j]), iw[q.z], q.z, half);
            if (b != curb) {
                if (curb >= 0) atomicAdd(&my[curb], acc);
                curb = b;
                acc = v;
            } else {
                acc += v;
            }
        }
        if (curb >= 0) atomicAdd(&my[curb], acc);
#pragma unroll
        for (int r = 0; r < RUN; r++) cur[r] = nxt[r];
        advance(p, B * RUN, rowlen, a.dims);
    }
    store_partial<WAVES>(a, hist);
}

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

// Adds each lane's (bin b, value v) to hist[b], summing equal-bin runs of
// neighbouring lanes first.  A run (segment) starts at lane 0, at a row start
// and wherever b changes; a segmented inclusive scan (Hillis-Steele, stopped
// once no segment is longer than the distance covered) leaves each segment's
// total in its last lane, which alone touches LDS.  Lanes with b < 0 carry
// nothing.  All 64 lanes must call it.
__device__ __forceinline__ void wave_segment_add(double *hist, int b, double v, bool row_start, int lane)
{
    const int bp = __shfl_up(b, 1, 64);
    const unsigned long long heads = __ballot(lane == 0 || row_start || b != bp);
    // first lane of this lane's segment: the highest head at or below it
    const int s = 63 - __clzll((long long)(heads & ((2ull << lane) - 1ull)));
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        if (!__any(lane - d >= s)) break;
        const double o = __shfl_up(v, d, 64);
        if (lane - d >= s) v += o;
    }
    const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
    if (tail && b >= 0) atomicAdd(&hist[b], v);
}

template <int NH>
__host__ __device__ constexpr size_t seg_smem_bytes(int dims, int nrbins)
{
    return sizeof(double) * ((size_t)NH * nrbins + dims) + sizeof(int) * (nrbins + 1);
}

// Coalesced sweep: wave w of a block takes chunks of 64*U consecutive modes
// (w, w + WAVES, ... within the block's row range); load u of a chunk is one
// 1 KB contiguous dwordx4 load across the wave (non-temporal when NT), the
// next chunk's loads are issued before the current one is binned.  Equal-bin
// runs across lanes are summed by wave_segment_add.  NH histogram copies in
// LDS (NH = WAVES: one per wavefront, so the sums are the same run to run;
// fewer copies when the bins do not fit, shared by several wavefronts).
template <int WAVES, int U, int NH, bool NT>
__global__ __launch_bounds__(64 * WAVES) void k_bin_power_seg(BinArgs a)
{
    static_assert(WAVES % NH == 0, "histogram copies must divide the wavefronts");
    extern __shared__ double smem[];
    double *hist = smem;
    double *iw = hist + (size_t)NH * a.nrbins;
    int *thr = (int *)(iw + a.dims);
    for (int t = threadIdx.x; t < NH * a.nrbins; t += 64 * WAVES) hist[t] = 0.0;
    for (int t = threadIdx.x; t < a.dims; t += 64 * WAVES) iw[t] = a.iw[t];
    for (int t = threadIdx.x; t <= a.nrbins; t += 64 * WAVES) thr[t] = a.thr[t];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rowlen = a.dims / 2 + 1, half = a.dims / 2;
    const long long r0 = a.rows * blockIdx.x / gridDim.x, r1 = a.rows * (blockIdx.x + 1) / gridDim.x;
    const long long m0 = r0 * rowlen, m1 = r1 * rowlen;
    constexpr int CH = 64 * U;
    const long long S = (long long)CH * WAVES;
    const Step s64 = step_of(64, rowlen, a.dims), sS = step_of(S, rowlen, a.dims);
    double *my = hist + (size_t)(wave % NH) * a.nrbins;
    long long c = m0 + (long long)wave * CH;
    Pos p = pos_of((long long)wave * CH + lane, r0, rowlen, a.dims);
    double2 cur[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long long mu = c + 64 * u + lane;
        cur[u] = mu < m1 ? kn_mem::ld<NT>(a.F + mu) : make_double2(0.0, 0.0);
    }
    for (; c < m1; c += S) {
        double2 nxt[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long mu = c + S + 64 * u + lane;
            nxt[u] = mu < m1 ? kn_mem::ld<NT>(a.F + mu) : make_double2(0.0, 0.0);
        }
        Pos q = p;
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (u) advance(q, s64, rowlen, a.dims);
            const long long mu = c + 64 * u + lane;
            int b = -1;
            double v = 0.0;
            if (mu < m1) {
                const int kx = kval(a.startslab + q.i, a.dims), ky = kval(q.j, a.dims);
                const int k2 = kx * kx + ky * ky + q.z * q.z;
                if (k2 != 0) {
                    b = find_bin(k2, a.half_bpu, thr, a.nrbins);
                    v = mode_weight(cur[u], __dmul_rn(iw[a.startslab + q.i], iw[q.j]), iw[q.z], q.z, half);
                }
            }
            wave_segment_add(my, b, v, q.z == 0, lane);
        }
#pragma unroll
        for (int u = 0; u < U; u++) cur[u] = nxt[u];
        advance(p, sS, rowlen, a.dims);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < a.nrbins; b += 64 * WAVES) {
        double s = hist[b];
        for (int w = 1; w < NH; w++) s += hist[(size_t)w * a.nrbins + b];
        a.partial[(size_t)blockIdx.x * a.nrbins + b] = s;
    }
}

// Slot of mode m (0 <= m < 64*RUN) of a wavefront's staging buffer.  Lanes
// write consecutive modes (m = 64u + l) and read back runs (m = RUN l + r).
// A 16-byte access pass serves 8 or 16 lanes; for the reads, the lanes of a
// pass differ in the bits of l above the run length, so those bits (folded
// so that both pass widths see a one-to-one map) are XOR-ed into the slot's
// low bits; the writes of a pass keep distinct low bits since the XOR term is
// constant across each aligned group of 8 lanes.  The slot stays inside m's
// aligned group of 2, 4 or 8 slots.
template <int RUN>
__device__ __forceinline__ int stage_slot(int m)
{
    static_assert(RUN == 2 || RUN == 4 || RUN == 8, "runs of 2, 4 or 8");
    if constexpr (RUN == 2)  // l bits l0..l3 = m1..m4: XOR l2 ^ l3 into m0
        return m ^ (((m >> 3) ^ (m >> 4)) & 1);
    else if constexpr (RUN == 4)  // l bits l0..l3 = m2..m5: XOR (l1 ^ l3, l2) into m0, m1
        return m ^ ((((m >> 3) ^ (m >> 5)) & 1) | (((m >> 4) & 1) << 1));
    else  // l bits l0..l3 = m3..m6: XOR (l0 ^ l3, l1, l2) into m0..m2
        return m ^ ((((m >> 3) ^ (m >> 6)) & 1) | (((m >> 4) & 1) << 1) | (((m >> 5) & 1) << 2));
}

// Coalesced loads, per-lane runs: wave w takes chunks of 64*RUN consecutive
// modes (w, w + WAVES, ... within the block's row range).  Load u of a chunk
// is one 1 KB contiguous dwordx4 load across the wave (non-temporal when NT);
// the chunk is passed through the wave's LDS staging buffer so that lane l
// holds modes 4l .. 4l+RUN-1 of it, and each lane bins its run as
// k_bin_power_runs_pf does (incremental bin search, equal-bin runs summed in
// registers, one LDS atomic per run).  The next chunk's loads are issued
// before the current one is staged.  NH histogram copies (NH = WAVES: one
// per wavefront, sums the same run to run).
template <int WAVES, int RUN, int NH, bool NT>
__host__ __device__ constexpr size_t lds_smem_bytes(int dims, int nrbins)
{
    return sizeof(double) * ((size_t)NH * nrbins + dims) + sizeof(int) * (nrbins + 1) +
           sizeof(double2) * (size_t)WAVES * 64 * RUN;
}

template <int WAVES, int RUN, int NH, bool NT, int PF = 1>
__global__ __launch_bounds__(64 * WAVES) void k_bin_power_lds(BinArgs a)
{
    static_assert(WAVES % NH == 0, "histogram copies must divide the wavefronts");
    extern __shared__ double smem[];
    double2 *stage_all = (double2 *)smem;  // WAVES * 64 * RUN modes, 16-byte aligned
    double *hist = (double *)(stage_all + (size_t)WAVES * 64 * RUN);
    double *iw = hist + (size_t)NH * a.nrbins;
    int *thr = (int *)(iw + a.dims);
    for (int t = threadIdx.x; t < NH * a.nrbins; t += 64 * WAVES) hist[t] = 0.0;
    for (int t = threadIdx.x; t < a.dims; t += 64 * WAVES) iw[t] = a.iw[t];
    for (int t = threadIdx.x; t <= a.nrbins; t += 64 * WAVES) thr[t] = a.thr[t];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double2 *stage = stage_all + (size_t)wave * 64 * RUN;
    const int rowlen = a.dims / 2 + 1, half = a.dims / 2;
    const long long r0 = a.rows * blockIdx.x / gridDim.x, r1 = a.rows * (blockIdx.x + 1) / gridDim.x;
    const long long m0 = r0 * rowlen, m1 = r1 * rowlen;
    constexpr int CH = 64 * RUN;
    const long long S = (long long)CH * WAVES;
    const Step sS = step_of(S, rowlen, a.dims);
    double *my = hist + (size_t)(wave % NH) * a.nrbins;
    long long c = m0 + (long long)wave * CH;
    Pos p = pos_of((long long)wave * CH + (long long)lane * RUN, r0, rowlen, a.dims);
    // ring of PF + 1 chunks: buf[0] is binned while buf[1..PF] are in flight
    double2 buf[PF + 1][RUN];
#pragma unroll
    for (int k = 0; k < PF; k++)
#pragma unroll
        for (int u = 0; u < RUN; u++) {
            const long long mu = c + k * S + 64 * u + lane;
            buf[k][u] = mu < m1 ? kn_mem::ld<NT>(a.F + mu) : make_double2(0.0, 0.0);
        }
    for (; c < m1; c += S) {
#pragma unroll
        for (int u = 0; u < RUN; u++) {
            const long long mu = c + PF * S + 64 * u + lane;
            buf[PF][u] = mu < m1 ? kn_mem::ld<NT>(a.F + mu) : make_double2(0.0, 0.0);
        }
        // transpose through LDS: lane l wrote modes l, 64+l, ...; reads RUN l .. RUN l + RUN - 1
#pragma unroll
        for (int u = 0; u < RUN; u++) stage[stage_slot<RUN>(64 * u + lane)] = buf[0][u];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double2 f[RUN];
#pragma unroll
        for (int r = 0; r < RUN; r++) f[r] = stage[stage_slot<RUN>(RUN * lane + r)];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const long long mrun = c + (long long)RUN * lane;
        Pos q = p;
        int curb = -1, b = -1, bnext = 0;
        double acc = 0.0, iwxy = 0.0;
#pragma unroll
        for (int r = 0; r < RUN; r++) {
            if (mrun + r >= m1) break;
            if (r) advance(q, 1, rowlen, a.dims);
            // the row's window product, at the run start and at a new row
            if (r == 0 || q.z == 0) iwxy = __dmul_rn(iw[a.startslab + q.i], iw[q.j]);
            const int kx = kval(a.startslab + q.i, a.dims), ky = kval(q.j, a.dims);
            const int k2 = kx * kx + ky * ky + q.z * q.z;
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
                if (curb >= 0) atomicAdd(&my[curb], acc);
                curb = b;
                acc = v;
            } else {
                acc += v;
            }
        }
        if (curb >= 0) atomicAdd(&my[curb], acc);
#pragma unroll
        for (int k = 0; k < PF; k++)
#pragma unroll
            for (int u = 0; u < RUN; u++) buf[k][u] = buf[k + 1][u];
        advance(p, sS, rowlen, a.dims);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < a.nrbins; b += 64 * WAVES) {
        double s = hist[b];
        for (int w = 1; w < NH; w++) s += hist[(size_t)w * a.nrbins + b];
        a.partial[(size_t)blockIdx.x * a.nrbins + b] = s;
    }
}

}  // namespace kn_bin

#endif
