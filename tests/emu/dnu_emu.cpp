// The library's per-call free-streaming kernels (kspace-neutrinos_amd/csrc/delta_nu.hip)
// built for the CPU SIMT emulation (include/hip/hip_runtime.h).  Test infrastructure only.
#include <hip/hip_runtime.h>
unsigned emu_tid, emu_bid, emu_bdim, emu_gdim;
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
