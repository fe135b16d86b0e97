// The library's per-call free-streaming kernels (kspace-neutrinos_amd/csrc/delta_nu.hip)
// built for the CPU SIMT emulation (include/hip/hip_runtime.h).  Test infrastructure only.
#include <hip/hip_runtime.h>
unsigned emu_tid, emu_bid, emu_bdim, emu_gdim, emu_bid_y, emu_gdim_y = 1;
EmuWaveSched *emu_wave;
int emu_cus = 2, emu_per_cu = 1;
double emu_dyn_lds[160 * 1024 / sizeof(double)];

// Stack switch (x86-64 System V): push callee-saved registers, swap stack
// pointers, pop the other context's registers and return into it.
asm(R"(
    .text
    .globl emu_ctx_switch
    .type emu_ctx_switch, @function
emu_ctx_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_ctx_switch, .-emu_ctx_switch
)");

// First code a lane fiber runs: the kernel body, then back to the scheduler for good.
extern "C" void emu_fiber_entry()
{
    EmuWaveSched &S = *emu_wave;
    S.body(S.arg);
    S.done[S.cur] = true;
    void *dead;
    emu_ctx_switch(&dead, S.main_sp);
}

#include "delta_nu.hip"

// ---- test hook: the two-pass fslength nodes against one full QAG per node ----
namespace {
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_fs_full_test(HTab H, int nfs, double x0, double xa,
                                                                       double light, double *fx, double *fy,
                                                                       int *status)
{
    __shared__ QagWs ws[kWavesPerBlock];
    const int wave = threadIdx.x >> 6;
    const int node = blockIdx.x * kWavesPerBlock + wave;
    if (node >= nfs) return;
    fs_node_wave(H, nfs, x0, xa, light, fx, fy, status, node, ws[wave]);
}
}  // namespace

// fslength at nfs nodes over [x0, xa] with the H(a) of hub(a, ctx), by the
// two passes the library uses (two_pass) and by one wavefront QAG per node
// (full); returns the number of nodes the first pass handed to the second.
extern "C" int emu_fs_two_pass_check(double (*hub)(double, void *), void *ctx, double x0, double xa, int nfs,
                                     double light, double *two_pass, double *full)
{
    upload_constants();
    const double w = 1.0 / 256.0;
    const int npanel = (int)ceil((xa - x0) / w) + 1;
    std::vector<double> coef;
    HTab H;
    h_table(x0, w, npanel, [hub, ctx](double a) { return hub(a, ctx); }, coef, H);
    H.coef = coef.data();
    std::vector<double> fx(nfs), fy(nfs), gx(nfs);
    std::vector<int> list(nfs + 1, 0), status(2, 0);
    hipLaunchKernelGGL(k_fs_first, dim3(fs_first_blocks(nfs)), dim3(256), 0, 0, H, nfs, x0, xa, light, fx.data(),
                       fy.data(), status.data(), list.data(), list.data() + nfs);
    const int listed = list[nfs];
    hipLaunchKernelGGL(k_fs_rest, dim3(kFsRestBlocks), dim3(64 * kWavesPerBlock), 0, 0, H, nfs, x0, xa, light,
                       fx.data(), fy.data(), status.data(), list.data(), list.data() + nfs);
    memcpy(two_pass, fy.data(), sizeof(double) * nfs);
    hipLaunchKernelGGL(k_fs_full_test, dim3(blocks_for(nfs)), dim3(64 * kWavesPerBlock), 0, 0, H, nfs, x0, xa, light,
                       gx.data(), full, status.data());
    return listed;
}

// ---- test hook: the wavefront error-order update against QUADPACK's serial one ----
namespace {
__global__ __launch_bounds__(64) void k_qag_sort_test(const double *errs, int nsteps, int limit, int *bad)
{
    __shared__ QagWs w1, w2;  // serial (lane 0) and wavefront
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        w1.e[0] = w2.e[0] = 1.0;
        w1.ord[0] = w2.ord[0] = 0;
    }
    int nrmax = 0, imax1 = 0, imax2 = __shfl(0, 0, 64);
    for (int t = 0; t < nsteps && t + 1 < limit; t++) {
        const int n = t + 1;  // intervals before this bisection
        if (lane == 0) {
            w1.e[imax1] = errs[2 * t];
            w1.e[n] = errs[2 * t + 1];
            w2.e[imax2] = errs[2 * t];
            w2.e[n] = errs[2 * t + 1];
            qag_sort(w1, n + 1, limit, nrmax, imax1);
        }
        imax2 = qag_sort_wave(w2, n + 1, limit);
        imax1 = __shfl(imax1, 0, 64);
        if (lane == 0) {
            int diff = imax1 != imax2;
            for (int q = 0; q <= n; q++) diff |= w1.ord[q] != w2.ord[q];
            bad[t] = diff;
        }
        __builtin_amdgcn_wave_barrier();
    }
}
}  // namespace

// Runs nsteps bisections whose two new error estimates are errs[2t], errs[2t+1]
// through qag_sort and qag_sort_wave; bad[t] = 1 where their lists differ.
extern "C" void emu_qag_sort_check(const double *errs, int nsteps, int limit, int *bad)
{
    hipLaunchKernelGGL(k_qag_sort_test, dim3(1), dim3(64), 0, 0, errs, nsteps, limit, bad);
}

// The device rule's error rescaling (dnu_device.h gk_rescale), for tests.
extern "C" double emu_gk_rescale(double err, double resabs, double resasc)
{
    return kn_dnu::gk_rescale(err, resabs, resasc);
}
