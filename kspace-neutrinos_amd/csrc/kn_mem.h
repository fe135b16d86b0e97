// kn_mem.h -- 16-byte loads and stores of one FP64 complex mode, with an
// optional non-temporal hint (global_load/store_dwordx4 ... nt), for the
// grid sweeps: each mode is touched once per sweep, so the hint keeps the
// stream from evicting the tables and partial sums the sweeps reuse.
#ifndef KN_MEM_H
#define KN_MEM_H

#include <hip/hip_runtime.h>

namespace kn_mem {

typedef double v2d __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ double2 ld(const double2 *p)
{
    if constexpr (NT) {
        const v2d v = __builtin_nontemporal_load((const v2d *)p);
        return make_double2(v.x, v.y);
    } else {
        return *p;
    }
}

template <bool NT>
__device__ __forceinline__ void st(double2 *p, double2 f)
{
    if constexpr (NT) {
        v2d v;
        v.x = f.x;
        v.y = f.y;
        __builtin_nontemporal_store(v, (v2d *)p);
    } else {
        *p = f;
    }
}

}  // namespace kn_mem

#endif
