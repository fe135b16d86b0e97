/* kn_testhost.c -- the N-body host services the library imports
 * (gadget_defines.h:11-32), as the reference's test suite provides them:
 * terminate/message/mymalloc/myfree follow gadget_defines_nompi.c:11-47 and
 * hubble_function is the test form of delta_tot_table_test.c:25-45
 * (matter + Lambda, no curvature, plus Omega_nu(a) and photons), built on the
 * library's own get_omega_nu/get_omegag as the reference test does.
 * Test infrastructure, loaded with RTLD_GLOBAL before the library. */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <math.h>

#include "kspace_neutrinos_hip.h"

static const _omega_nu *h_omnu;
static double h_Omega_nonu, h_OmegaLambda, h_Hubble;
static long long n_hubble_calls, n_messages;

void kn_test_init_hubble(const _omega_nu *omnu, double Omega0, double UnitTime_in_s)
{
    h_omnu = omnu;
    h_Omega_nonu = Omega0 - get_omega_nu(omnu, 1);
    h_OmegaLambda = 1 - Omega0;
    h_Hubble = 3.24077929e-18 * UnitTime_in_s;
}

double hubble_function(double a)
{
    double omega_tot;
    n_hubble_calls++;
    if (!h_omnu) terminate(1, "kn_test_init_hubble was not called before hubble_function\n");
    omega_tot = h_Omega_nonu / pow(a, 3) + h_OmegaLambda;
    omega_tot += get_omega_nu(h_omnu, a);
    omega_tot += get_omegag(h_omnu, a);
    return h_Hubble * sqrt(omega_tot);
}

/* Per-history form of the same H(a) for batched solves: ctx is a kn_test_hctx. */
typedef struct {
    const _omega_nu *omnu;
    double Omega_nonu, OmegaLambda, Hubble;
} kn_test_hctx;

void kn_test_hctx_init(kn_test_hctx *c, const _omega_nu *omnu, double Omega0, double UnitTime_in_s)
{
    c->omnu = omnu;
    c->Omega_nonu = Omega0 - get_omega_nu(omnu, 1);
    c->OmegaLambda = 1 - Omega0;
    c->Hubble = 3.24077929e-18 * UnitTime_in_s;
}

double kn_test_hubble_ctx(double a, void *ctx)
{
    const kn_test_hctx *c = (const kn_test_hctx *)ctx;
    double omega_tot = c->Omega_nonu / pow(a, 3) + c->OmegaLambda;
    omega_tot += get_omega_nu(c->omnu, a);
    omega_tot += get_omegag(c->omnu, a);
    return c->Hubble * sqrt(omega_tot);
}

long long kn_test_hubble_calls(void) { return n_hubble_calls; }
long long kn_test_messages(void) { return n_messages; }

void terminate(int ierr, const char *fmt, ...)
{
    va_list va;
    va_start(va, fmt);
    fprintf(stderr, "terminate(%d): ", ierr);
    vfprintf(stderr, fmt, va);
    va_end(va);
    fflush(stderr);
    exit(ierr);
}

void message(int ierr, const char *fmt, ...)
{
    n_messages++;
    if (getenv("KN_VERBOSE")) {
        va_list va;
        va_start(va, fmt);
        vprintf(fmt, va);
        va_end(va);
    }
    (void)ierr;
}

void *mymalloc_fullinfo(const char *string, size_t size, const char *func, const char *file, int line)
{
    (void)string; (void)func; (void)file; (void)line;
    return malloc(size);
}

void myfree_fullinfo(void *ptr, const char *func, const char *file, int line)
{
    (void)func; (void)file; (void)line;
    free(ptr);
}
