/* kn_hip.h -- helpers shared by the HIP sources (gfx950 only). */
#ifndef KN_HIP_H
#define KN_HIP_H

#include <hip/hip_runtime.h>

#include "kn_internal.h"

#define KN_CHECK(call)                                                                          \
    do {                                                                                        \
        hipError_t kn_e_ = (call);                                                              \
        if (kn_e_ != hipSuccess) {                                                              \
            kn_set_last_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(kn_e_), __FILE__, \
                              __LINE__);                                                        \
            kn_terminate(KN_ERR_HIP, "HIP error %s at %s:%d: %s\n", hipGetErrorString(kn_e_),   \
                         __FILE__, __LINE__, #call);                                            \
        }                                                                                       \
    } while (0)

/* KVAL of powerspectrum.c:27: FFT storage index -> signed wavenumber. */
__host__ __device__ static inline int kn_kval(int n, int dims) { return n <= dims / 2 ? n : n - dims; }

/* Is p device-accessible memory (device or managed allocation)? */
static inline int kn_is_device_ptr(const void *p)
{
    hipPointerAttribute_t at;
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

#endif
