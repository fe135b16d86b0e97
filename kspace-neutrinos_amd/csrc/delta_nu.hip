// delta_nu.hip -- placeholder until the integrator kernels land.
#include "kn_hip.h"
extern "C" void kn_dnu_solve(const _delta_tot_table *, double, const double *, int, const kn_species *, double *)
{
    kn_terminate(KN_ERR_ARG, "get_delta_nu: GPU integrator not built\n");
}
extern "C" void kn_dnu_release(const _delta_tot_table *) {}
