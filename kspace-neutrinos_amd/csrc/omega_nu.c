/* omega_nu.c -- neutrino and photon background densities (host only).
 *
 * Same API and algorithm as omega_nu_single.c:16-279.  These are O(1)
 * scalars per step; they feed the integrator's prefactors and the host's
 * hubble_function, so they stay on the host.  The Fermi-Dirac density table
 * is 200 points in log a, each a QAG-GK61 integral at relative error 1e-9
 * over q in [0, 500 kT], interpolated with a natural cubic spline.
 */
#include <math.h>
#include <string.h>

#include "kn_internal.h"

#define NRHOTAB 200
#define NU_SW 100 /* switch to the analytic series when kT/(a m) < 1/NU_SW */

static double tnu_over_tcmb(void) { return pow(4 / 11., 1 / 3.) * 1.00328; } /* omega_nu_single.h:17 */

/* (eV/c)^4 -> g cm^-3 for one species incl. antineutrinos (omega_nu_single.c:100-114) */
static double rho_conversion(void)
{
    const double chbar = 1. / (2 * M_PI * KN_LIGHTCGS * KN_HBAR);
    double conv = 4 * M_PI * 2;
    conv *= chbar * chbar * chbar;
    conv *= 1.60217646e-12 / KN_LIGHTCGS / KN_LIGHTCGS;
    return conv;
}

typedef struct { double am, kT; } fd_par;

static double fd_energy_integrand(double q, void *vp)
{
    const fd_par *p = (const fd_par *)vp;
    const double eps = sqrt(q * q + p->am * p->am);
    const double occ = 1. / (exp(q / p->kT) + 1);
    return q * q * eps * occ;
}

double kn_rho_conversion(void) { return rho_conversion(); }

double kn_kbtnu(double tcmb0) { return KN_BOLEVK * tnu_over_tcmb() * tcmb0; }

double kn_rhocrit(double HubbleParam)
{
    return (3 * KN_HUBBLE * HubbleParam * KN_HUBBLE * HubbleParam) / (8 * M_PI * KN_GRAVITY);
}

double kn_omegag0(double tcmb0, double rhocrit)
{
    return 4 * KN_STEFAN_BOLTZMANN / (KN_LIGHTCGS * KN_LIGHTCGS * KN_LIGHTCGS) * pow(tcmb0, 4) / rhocrit;
}

/* omega_nu_single.c:121-130: the table spans a0/1.2 to 1.2 NU_SW kT/m in a
 * (the series takes over above NU_SW kT/m); none for a nearly massless
 * species or an empty range. */
int kn_rho_table_range(double a0, double mnu, double kBtnu, double *lo, double *hi)
{
    *lo = log(a0) - log(1.2);
    *hi = log(NU_SW * kBtnu / mnu) + log(1.2);
    return !(mnu < 1e-6 * kBtnu || *hi < *lo);
}

void rho_nu_init(_rho_nu_single *t, double a0, const double mnu, const double HubbleParam, const double kBtnu)
{
    double lo, hi;
    int i;
    (void)HubbleParam;
    t->mnu = mnu;
    t->loga = t->rhonu = NULL;
    t->interp = t->acc = NULL;
    if (!kn_rho_table_range(a0, mnu, kBtnu, &lo, &hi)) return;
    t->loga = kn_malloc("rho_nu_table", 2 * NRHOTAB * sizeof(double));
    t->rhonu = t->loga + NRHOTAB;
    t->interp = kn_malloc("rho_nu_spline", NRHOTAB * sizeof(double));
    for (i = 0; i < NRHOTAB; i++) {
        fd_par par;
        double val, err;
        t->loga[i] = lo + i * (hi - lo) / (NRHOTAB - 1);
        par.am = mnu * exp(t->loga[i]);
        par.kT = kBtnu;
        if (kn_qag61(fd_energy_integrand, &par, 0, 500 * kBtnu, 0, 1e-9, KN_GSL_VAL, &val, &err))
            kn_terminate(2001, "quadrature failure in rho_nu_init\n");
        t->rhonu[i] = val / pow(exp(t->loga[i]), 4) * rho_conversion();
    }
    kn_cspline_init(t->loga, t->rhonu, NRHOTAB, (double *)t->interp);
}

static double rho_nonrel(double a, double kT, double am, double r2)
{
    /* zeta(3), zeta(5), zeta(7) series (omega_nu_single.c:156-160) */
    return am * (kT * kT * kT) / (a * a * a * a) *
           (1.5 * 1.202056903159594 + r2 * 45. / 4. * 1.0369277551433704 +
            2835. / 32. * r2 * r2 * 1.0083492773819229 + 80325 / 32. * r2 * r2 * r2 * 1.0020083928260826) *
           rho_conversion();
}

static double rho_rel(double a, double kT) { return 7 * pow(M_PI * kT / a, 4) / 120. * rho_conversion(); }

double rho_nu(_rho_nu_single *t, const double a, const double kT)
{
    const double am = a * t->mnu;
    const double r2 = kT * kT / am / am;
    double la;
    if (NU_SW * NU_SW * r2 < 1) return rho_nonrel(a, kT, am, r2);
    if (am < 1e-6 * kT) return rho_rel(a, kT);
    la = log(a);
    if (!t->loga || la < t->loga[0]) return am < 1e-4 * kT ? rho_rel(a, kT) : rho_nonrel(a, kT, am, r2);
    {
        int edom = 0;
        const double v = kn_cspline_eval(t->loga, t->rhonu, (const double *)t->interp, NRHOTAB, la, &edom);
        if (edom) kn_terminate(2001, "rho_nu: log a = %g outside the density table\n", la);
        return v;
    }
}

/* species of equal mass (within FLOAT_ACC) are folded into the first one
 * (omega_nu_single.c:27-40) */
void kn_nu_degeneracies(const double MNu[], int deg[])
{
    int i;
    for (i = 0; i < NUSPECIES; i++) {
        int j;
        deg[i] = 0;
        for (j = 0; j < i; j++)
            if (fabs(MNu[i] - MNu[j]) < KN_FLOAT_ACC) {
                deg[j] += 1;
                break;
            }
        if (j == i) deg[i] = 1;
    }
}

void init_omega_nu(_omega_nu *om, const double MNu[], const double a0, const double HubbleParam,
                   const double tcmb0)
{
    int i;
    memset(&om->hybnu, 0, sizeof om->hybnu);
    om->tcmb0 = tcmb0;
    om->kBtnu = kn_kbtnu(tcmb0);
    om->rhocrit = kn_rhocrit(HubbleParam);
    kn_nu_degeneracies(MNu, om->nu_degeneracies);
    for (i = 0; i < NUSPECIES; i++) {
        if (om->nu_degeneracies[i]) {
            om->RhoNuTab[i] = kn_malloc("RhoNuTab", sizeof(_rho_nu_single));
            rho_nu_init(om->RhoNuTab[i], a0, MNu[i], HubbleParam, om->kBtnu);
        } else {
            om->RhoNuTab[i] = NULL;
        }
    }
}

double get_omega_nu(const _omega_nu *om, const double a)
{
    double rho = 0;
    int i;
    for (i = 0; i < NUSPECIES; i++)
        if (om->nu_degeneracies[i] > 0) rho += om->nu_degeneracies[i] * rho_nu(om->RhoNuTab[i], a, om->kBtnu);
    return rho / om->rhocrit;
}

double particle_nu_fraction(const _hybrid_nu *h, const double a, int i)
{
    if (!h->enabled) return 0;
    return a > h->nu_crit_time ? h->nufrac_low[i] : 0;
}

double get_omega_nu_nopart(const _omega_nu *om, const double a)
{
    const double all = get_omega_nu(om, a);
    const double part = get_omega_nu(om, 1) * particle_nu_fraction(&om->hybnu, a, 0) / (a * a * a);
    return all - part;
}

double get_omegag(const _omega_nu *om, const double a)
{
    return kn_omegag0(om->tcmb0, om->rhocrit) / pow(a, 4);
}

double omega_nu_single(const _omega_nu *om, const double a, int i)
{
    double onu, opart;
    if (om->nu_degeneracies[i] == 0) {
        int j;
        for (j = i; j >= 0; j--)
            if (om->nu_degeneracies[j]) {
                i = j;
                break;
            }
    }
    onu = rho_nu(om->RhoNuTab[i], a, om->kBtnu) / om->rhocrit;
    opart = rho_nu(om->RhoNuTab[i], 1, om->kBtnu) / om->rhocrit;
    opart *= particle_nu_fraction(&om->hybnu, a, i) / (a * a * a);
    return onu - opart;
}

static double fd_number_integrand(double x, void *p)
{
    (void)p;
    return x * x / (exp(x) + 1);
}

double nufrac_low(const double qc)
{
    double v, err;
    if (kn_qag61(fd_number_integrand, NULL, 0, qc, 0, 1e-6, 100, &v, &err))
        kn_terminate(2001, "quadrature failure in nufrac_low\n");
    return v / (1.5 * KN_ZETA3);
}

void init_hybrid_nu(_hybrid_nu *h, const double mnu[], const double vcrit, const double light,
                    const double nu_crit_time, const double kBtnu)
{
    int i;
    h->enabled = 1;
    h->nu_crit_time = nu_crit_time;
    h->vcrit = vcrit / light;
    for (i = 0; i < NUSPECIES; i++) h->nufrac_low[i] = nufrac_low(mnu[i] * vcrit / light / kBtnu);
}
