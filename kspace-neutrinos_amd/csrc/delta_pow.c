/* delta_pow.c -- the delta_nu/delta_cdm ratio table (host side).
 *
 * Same API and semantics as delta_pow.c:6-43 of the reference: init stores
 * pointers to the caller's log k and ratio arrays (no copy) and a
 * normalisation; get_dnudcdm_powerspec clamps log k into the table (values
 * below logkk[0] - log 2 are reported on stderr, as the reference does) and
 * interpolates linearly.  The per-mode use inside add_nu_power_to_rhogrid runs
 * on the device (rescale.hip); this host function serves callers that query
 * single wavenumbers.
 */
#include <math.h>
#include <stdio.h>

#include "kn_internal.h"

void init_delta_pow(_delta_pow *d_pow, double logkk[], double delta_ratio[], int nbins, double norm)
{
    d_pow->logkk = logkk;
    d_pow->delta_ratio = delta_ratio;
    d_pow->nbins = nbins;
    d_pow->norm = norm;
    d_pow->spline = NULL;
    d_pow->acc = NULL;
}

double get_dnudcdm_powerspec(_delta_pow *d_pow, double kk)
{
    const double lo = d_pow->logkk[0];
    int edom = 0;
    double v;
    if (kk < lo && kk > lo - log(2)) kk = lo;
    if (kk < lo - log(2)) {
        fprintf(stderr, "get_dnudcdm_powerspec: log k = %g below the table (from %g); clamped\n", kk, lo);
        kk = lo;
    }
    if (kk > d_pow->logkk[d_pow->nbins - 1]) kk = d_pow->logkk[d_pow->nbins - 1];
    v = kn_linear_eval(d_pow->logkk, d_pow->delta_ratio, d_pow->nbins, kk, &edom);
    if (edom) kn_terminate(2001, "get_dnudcdm_powerspec: log k = %g outside the table\n", kk);
    return d_pow->norm * v;
}

void free_d_pow(_delta_pow *d_pow)
{
    d_pow->spline = NULL;
    d_pow->acc = NULL;
}
