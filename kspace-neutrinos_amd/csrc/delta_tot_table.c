/* delta_tot_table.c -- integrator state and the get_delta_nu entry points.
 *
 * Same API, state layout and control flow as delta_tot_table.c:25-250 and
 * :375-617 of the reference: the history block scalefact[namax] followed by
 * delta_tot[k][namax] rows, the lazy init from the CAMB transfer table, the
 * per-step update that appends an estimated row, integrates, then keeps or
 * drops the row (kept when a >= a_prev + 0.009), the NaN abort (2004) and the
 * clamp of negative delta_nu to 0.  The free-streaming integral itself
 * (get_delta_nu, delta_tot_table.c:507-611) runs on the GPU through
 * kn_dnu_solve (delta_nu.hip); this file prepares the per-species parameters
 * (mass, hybrid cut, tolerance) exactly as the reference derives them.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kn_internal.h"

void allocate_delta_tot_table(_delta_tot_table *d, const int nk_in, const double TimeTransfer, const double TimeMax,
                              const double Omega0, const _omega_nu *const omnu, const double UnitTime_in_s,
                              const double UnitLength_in_cm, int debug)
{
    int k;
    d->nk_allocated = nk_in;
    d->nk = nk_in;
    d->TimeTransfer = TimeTransfer;
    d->namax = (int)ceil(100 * (TimeMax - TimeTransfer)) + 2;
    d->ia = 0;
    d->delta_tot_init_done = 0;
    d->ThisTask = 0;
    d->delta_tot = kn_malloc("kspace_delta_tot", (size_t)nk_in * sizeof(double *));
    d->scalefact = kn_malloc("kspace_scalefact", (size_t)d->namax * (size_t)(nk_in + 1) * sizeof(double));
    memset(d->scalefact, 0, (size_t)d->namax * (size_t)(nk_in + 1) * sizeof(double));
    for (k = 0; k < nk_in; k++) d->delta_tot[k] = d->scalefact + (size_t)d->namax * (size_t)(k + 1);
    d->delta_nu_init = kn_malloc("kspace_delta_nu_init", 3 * (size_t)nk_in * sizeof(double));
    memset(d->delta_nu_init, 0, 3 * (size_t)nk_in * sizeof(double));
    d->delta_nu_last = d->delta_nu_init + nk_in;
    d->wavenum = d->delta_nu_init + 2 * nk_in;
    d->omnu = omnu;
    d->light = KN_LIGHTCGS * UnitTime_in_s / UnitLength_in_cm;
    d->delta_nu_prefac = 1.5 * Omega0 * KN_HUBBLE * KN_HUBBLE * pow(UnitTime_in_s, 2) / d->light;
    d->Omeganonu = Omega0 - get_omega_nu(omnu, 1);
    d->debug = debug;
}

void kn_dnu_release(const _delta_tot_table *d);

void free_delta_tot_table(_delta_tot_table *d)
{
    kn_dnu_release(d);
    kn_free(d->delta_tot);
    kn_free(d->scalefact);
    kn_free(d->delta_nu_init);
}

double get_delta_tot(const double dnu, const double dcdm, const double OmegaNua3, const double Omeganonu,
                     const double Omeganu1, const double partnu)
{
    const double fcdm = 1 - OmegaNua3 / (Omeganonu + Omeganu1);
    return fcdm * (dcdm + dnu * OmegaNua3 / (Omeganonu + Omeganu1 * partnu));
}

/* ---- special functions (delta_tot_table.c:417-463) ---- */

static double sph_j0(double x)
{
    /* sin(x)/x; Taylor series to x^12 below |x| = 0.5 */
    if (fabs(x) < 0.5) {
        const double y = x * x;
        return 1.0 + y * (-1.0 / 6.0 + y * (1.0 / 120.0 + y * (-1.0 / 5040.0 + y * (1.0 / 362880.0 +
                         y * (-1.0 / 39916800.0 + y * (1.0 / 6227020800.0))))));
    }
    return sin(x) / x;
}

double specialJ(const double x, const double qc, const double nufrac_low_v)
{
    if (qc > 0) {
        /* Fermi-Dirac transform truncated below qc: 19-term series */
        double s = 0;
        int n;
        for (n = 1; n < 20; n++) {
            const double ii = (n * n + n * n * n * qc + n * qc * x * x - x * x) * qc * sph_j0(qc * x) +
                              (2 * n + n * n * qc + qc * x * x) * cos(qc * x);
            s += -1 * pow((-1), n) * exp(-n * qc) / (n * n + x * x) / (n * n + x * x) * ii;
        }
        return s / (1.5 * KN_ZETA3 * (1 - nufrac_low_v));
    }
    if (x <= 0.) return 1.;
    {
        const double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
        return (1. + 0.0168 * x2 + 0.0407 * x4) / (1. + 2.1734 * x2 + 1.6787 * exp(4.1811 * log(x)) + 0.1467 * x8);
    }
}

/* ---- free-streaming length (delta_tot_table.c:378-407) ---- */

static double fsl_integrand(double loga, void *p)
{
    const double a = exp(loga);
    (void)p;
    return 1. / a / (a * kn_hubble(a));
}

double fslength(const double logai, const double logaf, const double light)
{
    double v, e;
    if (logai >= logaf) return 0;
    if (kn_qag61(fsl_integrand, NULL, logai, logaf, 0, 1e-6, KN_GSL_VAL, &v, &e))
        kn_terminate(2001, "GSL_ERROR-equivalent: QAG failed in fslength(%g, %g)\n", logai, logaf);
    return light * v;
}

double kn_deriv_prefac(const _delta_tot_table *d)
{
    return d->TimeTransfer * (kn_hubble(d->TimeTransfer) / d->light) * d->TimeTransfer;
}

/* ---- species parameters as get_delta_nu derives them (delta_tot_table.c:513-551) ---- */

void kn_species_for(const _delta_tot_table *d, double a, double mnu, kn_species *s)
{
    const double partnu = particle_nu_fraction(&d->omnu->hybnu, a, 0);
    s->mnubykT = mnu / d->omnu->kBtnu;
    s->qc = 0;
    s->relerr = 1e-6;
    s->ic_only = 0;
    if (partnu > 0) {
        if (1 - partnu < 1e-3) s->ic_only = 1;
        s->qc = d->omnu->hybnu.vcrit * s->mnubykT;
        s->relerr /= (1. + 1e-5 - particle_nu_fraction(&d->omnu->hybnu, a, 0));
    }
    if (!(d->ia > 1 && s->mnubykT > 0)) s->ic_only = 1;
}

void get_delta_nu(const _delta_tot_table *const d, const double a, const double wavenum[], double delta_nu_curr[],
                  const double mnu)
{
    kn_species s;
    kn_species_for(d, a, mnu, &s);
    if (d->debug)
        kn_message(0, "get_delta_nu at a=%g over %d stored steps (k[0]=%g, last delta_tot[0]=%g, m_nu=%g)\n", a, d->ia,
                   wavenum[0], d->delta_tot[0][d->ia - 1], mnu);
    kn_dnu_solve(d, a, wavenum, 1, &s, delta_nu_curr);
}

void get_delta_nu_combined(const _delta_tot_table *const d, const double a, const double wavenum[],
                           double delta_nu_curr[])
{
    const double Omega_nu_tot = get_omega_nu_nopart(d->omnu, a);
    kn_species sp[NUSPECIES];
    double om[NUSPECIES];
    int idx[NUSPECIES], ns = 0, i, k;
    double *single;
    for (i = 0; i < NUSPECIES; i++) {
        if (d->omnu->nu_degeneracies[i] > 0) {
            om[ns] = d->omnu->nu_degeneracies[i] * omega_nu_single(d->omnu, a, i);
            kn_species_for(d, a, d->omnu->RhoNuTab[i]->mnu, &sp[ns]);
            idx[ns++] = i;
        }
    }
    memset(delta_nu_curr, 0, sizeof(double) * (size_t)d->nk);
    if (!ns) return;
    single = kn_malloc("delta_nu_single", sizeof(double) * (size_t)d->nk * (size_t)ns);
    kn_dnu_solve(d, a, wavenum, ns, sp, single);
    for (i = 0; i < ns; i++)
        for (k = 0; k < d->nk; k++) delta_nu_curr[k] += single[(size_t)i * d->nk + k] * om[i] / Omega_nu_tot;
    (void)idx;
    kn_free(single);
}

void update_delta_tot(_delta_tot_table *const d, const double a, const double delta_cdm_curr[],
                      const double delta_nu_curr[], const int overwrite)
{
    const double OmegaNua3 = get_omega_nu_nopart(d->omnu, a) * pow(a, 3);
    const double OmegaNu1 = get_omega_nu(d->omnu, 1);
    const double partnu = particle_nu_fraction(&d->omnu->hybnu, a, 0);
    int k;
    if (!overwrite) d->ia++;
    if (d->ia > d->namax) kn_terminate(KN_ERR_RANGE, "delta_tot table full: %d rows of %d\n", d->ia, d->namax);
    d->scalefact[d->ia - 1] = log(a);
    for (k = 0; k < d->nk; k++)
        d->delta_tot[k][d->ia - 1] =
            get_delta_tot(delta_nu_curr[k], delta_cdm_curr[k], OmegaNua3, d->Omeganonu, OmegaNu1, partnu);
}

void delta_tot_init(_delta_tot_table *const d, const int nk_in, const double wavenum[], const double delta_cdm_curr[],
                    const _transfer_init_table *const t, const double Time)
{
    int k;
    double *c;
    /* drop resumed rows at or after the current time (delta_tot_table.c:83-90) */
    for (k = 0; k < d->ia; k++) {
        if (log(Time) <= d->scalefact[k]) {
            d->ia = k;
            break;
        }
    }
    if (Time > d->TimeTransfer + 0.01 && d->ia == 0)
        kn_terminate(2023, "resume: the delta_tot table is empty although a resume state was expected\n");
    if (Time < d->TimeTransfer - 1e-5)
        kn_terminate(2024, "delta_tot_init at a=%g, before the transfer-function time %g\n", Time, d->TimeTransfer);
    if (nk_in > d->nk_allocated)
        kn_terminate(2011, "%d k bins given, the table holds only %d\n", nk_in, d->nk_allocated);
    d->nk = nk_in;
    {
        const double OmegaNua3 = get_omega_nu_nopart(d->omnu, d->TimeTransfer) * pow(d->TimeTransfer, 3);
        const double OmegaNu1 = get_omega_nu(d->omnu, 1);
        c = kn_malloc("transfer_spline", sizeof(double) * (size_t)t->NPowerTable);
        kn_cspline_init(t->logk, t->T_nu, t->NPowerTable, c);
        if (log(wavenum[d->nk - 1]) > t->logk[t->NPowerTable - 1])
            kn_terminate(2, "k = %g lies beyond the transfer table, which ends at k = %g\n", wavenum[d->nk - 1],
                         exp(t->logk[t->NPowerTable - 1]));
        for (k = 0; k < d->nk; k++) {
            int edom = 0;
            const double T = kn_cspline_eval(t->logk, t->T_nu, c, t->NPowerTable, log(wavenum[k]), &edom);
            const double OmegaMa = d->Omeganonu + OmegaNua3;
            if (edom) kn_terminate(2001, "transfer table interpolation at k = %g out of range\n", wavenum[k]);
            if (d->ia == 0) {
                const double partnu = particle_nu_fraction(&d->omnu->hybnu, d->TimeTransfer, 0);
                d->delta_tot[k][0] =
                    get_delta_tot(delta_cdm_curr[k] * T, delta_cdm_curr[k], OmegaNua3, d->Omeganonu, OmegaNu1, partnu);
            }
            d->delta_nu_init[k] = d->delta_tot[k][0] * OmegaMa / (OmegaMa - OmegaNua3 + T * OmegaNua3) * fabs(T);
            d->wavenum[k] = wavenum[k];
        }
        kn_free(c);
    }
    if (d->ia == 0) {
        d->scalefact[0] = log(d->TimeTransfer);
        d->ia = 1;
    }
    get_delta_nu_combined(d, exp(d->scalefact[d->ia - 1]), wavenum, d->delta_nu_last);
    d->delta_tot_init_done = 1;
}

void get_delta_nu_update(_delta_tot_table *const d, const double a, const int nk_in, const double keff[],
                         const double delta_cdm_curr[], double delta_nu_curr[], _transfer_init_table *t)
{
    int k;
    if (!d->delta_tot_init_done) delta_tot_init(d, nk_in, keff, delta_cdm_curr, t, a);
    if (!d->delta_tot_init_done) kn_terminate(2001, "delta_tot_init has not run on this table\n");
    if (nk_in != d->nk) kn_terminate(2002, "%d k bins given, the table stores %d\n", nk_in, d->nk);
    if (d->nk < 2) kn_terminate(2003, "only %d k bins: too few to interpolate\n", d->nk);
    if (log(a) - d->scalefact[d->ia - 1] < KN_FLOAT_ACC) {
        for (k = 0; k < d->nk; k++) delta_nu_curr[k] = d->delta_nu_last[k];
        return;
    }
    /* provisional row from the last delta_nu, then integrate */
    update_delta_tot(d, a, delta_cdm_curr, d->delta_nu_last, 0);
    get_delta_nu_combined(d, a, keff, delta_nu_curr);
    for (k = 0; k < d->nk; k++) d->delta_nu_last[k] = delta_nu_curr[k];
    if (a >= exp(d->scalefact[d->ia - 2]) + 0.009)
        update_delta_tot(d, a, delta_cdm_curr, delta_nu_curr, 1);
    else
        d->ia--;
    for (k = 0; k < d->nk; k++) {
        if (isnan(delta_nu_curr[k]))
            kn_terminate(2004, "NaN in the neutrino update: delta_nu=%g at k bin %d (delta_cdm=%g, k=%g)\n", delta_nu_curr[k], k,
                         delta_cdm_curr[k], keff[k]);
        if (delta_nu_curr[k] < 0) delta_nu_curr[k] = 0;
    }
}

/* ---- state files (delta_tot_table.c:254-374 of the reference): the same text
 * formats, so a run resumes from files the reference wrote and the reference
 * resumes from files written here.  One line per stored step:
 * "# <log a> <delta_tot[0]> ... <delta_tot[nk-1]>" (%le, 7 significant digits). */

static const char *state_path(const char *f) { return f ? f : "delta_tot_nu.txt"; }

void read_all_nu_state(_delta_tot_table *const d, char *savefile)
{
    const char *path = state_path(savefile);
    FILE *fd = fopen(path, "r");
    int row;
    if (!fd) return; /* nothing to resume */
    for (row = 0; row < d->namax; row++) {
        double la;
        int k;
        if (fscanf(fd, "# %lg ", &la) != 1) break;
        d->scalefact[row] = la;
        for (k = 0; k < d->nk; k++) {
            if (fscanf(fd, "%lg ", &d->delta_tot[k][row]) == 1) continue;
            /* a short first line sets nk; a short later line is an error */
            if (row != 0)
                kn_terminate(2006, "%d k values expected, %d read from %s at a=%g\n", d->nk, k, path,
                             exp(d->scalefact[row]));
            d->nk = k;
            break;
        }
    }
    /* the stored history must start at the transfer function's time */
    if (fabs(d->scalefact[0] - log(d->TimeTransfer)) > 1e-4)
        kn_terminate(2007, "%s starts with a=%g, transfer function is at a=%g\n", path, exp(d->scalefact[0]),
                     d->TimeTransfer);
    if (row > 0) d->ia = row;
    if (d->debug) kn_message(1, "%d stored steps read from %s\n", row, path);
    fclose(fd);
}

void save_delta_tot(const _delta_tot_table *const d, const int iia, char *savefile)
{
    const char *path = state_path(savefile);
    FILE *fd = fopen(path, "a");
    int k;
    if (!fd) kn_terminate(2012, "cannot write %s\n", path);
    fprintf(fd, "# %le ", d->scalefact[iia]);
    for (k = 0; k < d->nk; k++) fprintf(fd, "%le ", d->delta_tot[k][iia]);
    fprintf(fd, "\n");
    fclose(fd);
}

void save_all_nu_state(const _delta_tot_table *const d, char *savefile)
{
    const char *path = state_path(savefile);
    FILE *old = fopen(path, "r");
    int row;
    if (old) { /* keep the previous file as <name>.bak */
        const size_t n = strlen(path);
        char *bak = kn_malloc("kspace_state_bak", n + 5);
        fclose(old);
        memcpy(bak, path, n);
        memcpy(bak + n, ".bak", 5);
        rename(path, bak);
        kn_free(bak);
    }
    for (row = 0; row < d->ia; row++) save_delta_tot(d, row, (char *)path);
}

int save_nu_power(const _delta_tot_table *const d, const double Time, const int snapnum, const char *OutputDir)
{
    char path[1000];
    FILE *fd;
    int k;
    snprintf(path, sizeof path, "%s/powerspec_nu_%03d.txt", OutputDir, snapnum);
    if (!(fd = fopen(path, "w"))) {
        fprintf(stderr, "can't open file `%s` for writing\n", path);
        return -1;
    }
    fprintf(fd, "# k P_nu(k)\n");
    fprintf(fd, "# a = %g\n", Time);
    fprintf(fd, "# nbins = %d\n", d->nk);
    for (k = 0; k < d->nk; k++) fprintf(fd, "%g %g\n", d->wavenum[k], d->delta_nu_last[k] * d->delta_nu_last[k]);
    fclose(fd);
    return 0;
}
